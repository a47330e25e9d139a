"""The benchmark configurations at their full sizes on the GPU.

Config 2: a level-4 shell under the pressure ramp of config 1 (0 -> 2 p_cr,
mass damping 0.05, contact stiffness 100 E1 h, mu = 0.3, cells of 2h, dt =
0.5 x CFL): FP32 split forces against FP64 at checkpoints of the same run,
and the FP32 trajectory against the FP64 one.
Config 3: a level-6 shell (40 962 nodes, 81 920 elements), h = 0.005,
E2/E1 = 0.3, mu = 0.5, deflated to 30 % of its volume in FP32 split
precision with the reference's adaptive dt: at snapshots through the dense
self-contact, the sorted candidate pairs bit for bit against the oracle
rebuilt from the same positions, and the forces (FP64 and FP32) against the
oracle's.
Config 4: the 512-shell sweep (205 / 205 / 102 shells of levels 3 / 4 / 5)
under pressure ramps to 1.5-2.5 p_cr in one batch: one shell per level
against its oracle run, every shell finite and within its device lists.
Config 5: the learned warm start on 1024 unseen level-4 shells.
Printed figures are the measured errors (reported in the design notes)."""
import os
import time

import numpy as np
import pytest

import oracle
import shell_amd as S
from shell_amd import warmstart as W
from shell_amd.workloads import batch_sweep, radial_imperfection

pytestmark = pytest.mark.gpu

E1, NU = 1.0, 0.3


def say(*a):
    print(*a, flush=True)


def relerr(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


def pcr(h):
    return 2 * E1 * h * h / np.sqrt(3 * (1 - NU ** 2))


def _checker(X, T, prm, u, v, t):
    """a fresh context at a state: its own pair lists, no friction history,
    so evaluating forces there leaves the runs alone"""
    c = S.ShellSystem(X, T, **prm)
    c.set_state(u, v, np.zeros_like(u), t)
    c.apply_initial_conditions()
    return c


def test_config2_fp32_against_fp64_l4():
    """both runs take the reference's error-controlled dt (at most 0.5 x CFL):
    at a fixed dt the snap-through after buckling (near 0.35 p_cr for this
    orthotropic shell) drives one sheet through the other within a few
    hundred steps and the contact rows overflow, in both precisions; with
    the adaptive dt both runs go through the collapse into dense contact"""
    h = 0.01
    X, T = S.mesh_icosphere(4)
    X = radial_imperfection(X, 0.01 * h, seed=2)
    base = dict(E1=E1, E2=0.5, nu=NU, G12=0.35, thickness=h, rho=1.0, damping_alpha=0.05, contact=1,
                contact_stiffness=100 * E1 * h, mu_s=0.3, mu_d=0.3, cell_size_mult=4.0, cell_margin=0.1,
                rebuild_trigger=1)
    probe = S.ShellSystem(X, T, **base)
    dt = 1.0 / probe.cfl_omega_max(60)
    probe.close()
    steps, chunk = 50000, 5000
    prm = dict(base, dt=dt, dt_max=dt, pressure_final=-2 * pcr(h), pressure_ramp_time=steps * dt, adapt_dt_every=1,
               newmark_beta=0.0)
    g64 = S.ShellSystem(X, T, **prm)
    g32 = S.ShellSystem(X, T, **prm)
    z = np.zeros(X.size)
    for g, prec in ((g64, S.FP64), (g32, S.FP32_SPLIT)):
        g.set_state(z)
        g.apply_initial_conditions(precision=prec)
    V0 = g64.advance(0)["volume"]
    say(f"case config 2: L4, dt_max {dt:.4g}, pressure ramp to 2 p_cr over {steps} x dt_max, adaptive dt")
    t0 = time.perf_counter()
    rows = []
    for k in range(steps // chunk):
        r64 = g64.advance(chunk)
        r32 = g32.advance(chunk, precision=S.FP32_SPLIT)
        u64, v64, _, t = g64.get_state()
        u32, _, _, t32 = g32.get_state()
        c = _checker(X, T, prm, u64, v64, t)
        f64 = c.forces(u64, v64, t)[0]
        f32 = c.forces(u64, v64, t, precision=S.FP32_SPLIT)[0]
        c.close()
        ferr = np.linalg.norm(f32 - f64) / max(np.linalg.norm(f64), r64["fext_norm"])
        same_t = abs(t32 - t) <= 1e-6 * t
        terr = relerr(u32, u64) if same_t else np.nan
        rows.append((r64["volume"] / V0, ferr, terr, r32["volume"] / V0))
        say(f"  step {r64['n_steps']} t {t:.1f} / {t32:.1f}: force |f32 - f64| / max(|f64|, |f_ext|) {ferr:.2e}, "
            f"|u32 - u64| / |u64| {terr:.2e}, V/V0 {rows[-1][0]:.4f} / {rows[-1][3]:.4f}, contacts "
            f"{r64['n_contacts']} / {r32['n_contacts']}, rejected {r64['n_rejected']} / {r32['n_rejected']}, "
            f"wall {time.perf_counter() - t0:.1f} s")
        assert r64["problem"] < S.STOP_SIM and r32["problem"] < S.STOP_SIM
        assert r64["capacity_err"] == 0 and r32["capacity_err"] == 0
    assert r32["n_steps_fp32"] == steps and r64["n_steps"] == steps
    # forces at one state: the FP32 element path at its rounding level
    # (2e-5 of |f_ext| measured in the folded states of the collapse)
    assert max(r[1] for r in rows) < 1e-4
    # trajectories before the buckle: same times, FP32 within 1e-4
    pre = [r for r in rows if r[0] > 0.97]
    assert pre and all(r[2] < 1e-4 for r in pre)
    # both collapse into dense contact
    assert rows[-1][0] < 0.1 and rows[-1][3] < 0.1
    g64.close()
    g32.close()


def _oracle_at(X, T, prm, u, v, t):
    o = oracle.Oracle(X, T, **prm)
    o.set_state(u, v, np.zeros_like(u), t)
    o.init_dynamics()
    return o


def test_config3_l6_deflation_pairs_bitexact():
    h = 0.005
    X, T = S.mesh_icosphere(6)
    X = radial_imperfection(X, 0.01 * h, seed=6)
    base = dict(E1=E1, E2=0.3, nu=NU, G12=0.35, thickness=h, rho=1.0, damping_alpha=0.05, bulk_modulus=0.05,
                volume_final_ratio=0.3, volume_ramp_time=150.0, volume_at_prediction=1, contact=1, mu_s=0.5,
                mu_d=0.5, cell_size_mult=0.6, cell_margin=0.1, rebuild_trigger=1, newmark_beta=0.0)
    probe = S.ShellSystem(X, T, **base)
    dt = 1.0 / probe.cfl_omega_max(60)
    probe.close()
    prm = dict(base, dt=dt, dt_max=dt, adapt_dt_every=1)
    g = S.ShellSystem(X, T, **prm)
    g.set_state(np.zeros(X.size))
    g.apply_initial_conditions(precision=S.FP32_SPLIT)
    V0 = g.advance(0)["volume"]
    say(f"case config 3: L6 ({len(X)} nodes, {len(T)} elements), h {h}, deflation to 30 % over 150 time units")
    t0 = time.perf_counter()
    checked = 0
    for k in range(10):
        r = g.advance(2000, precision=S.FP32_SPLIT)
        ratio = r["volume"] / V0
        say(f"  steps {r['n_steps']} t {r['t']:.1f} V/V0 {ratio:.4f} pairs {r['n_pairs']} contacts {r['n_contacts']} "
            f"problem {r['problem']} wall {time.perf_counter() - t0:.1f} s")
        assert r["problem"] < S.STOP_SIM and r["capacity_err"] == 0
        if k not in (5, 7, 8, 9):
            continue
        # snapshot: a fresh GPU context and the oracle at the same positions
        u, v, _, t = g.get_state()
        fresh = _checker(X, T, prm, u, v, t)
        ng = fresh.compute_shell_shell_contact_candidates(force=True)
        e1g, e2g = fresh.pairs()
        o = _oracle_at(X, T, prm, u, v, t)
        no = o.rebuild(force=True)
        e1o, e2o = o.pairs()
        say(f"    snapshot: pairs GPU {ng} oracle {no}")
        assert ng == no
        assert np.array_equal(e1g, e1o) and np.array_equal(e2g, e2o)
        fo = o.forces(u, v, t)[0]
        f64 = fresh.forces(u, v, t)[0]
        f32 = fresh.forces(u, v, t, precision=S.FP32_SPLIT)[0]
        say(f"    forces: FP64 {relerr(f64, fo):.2e}, FP32 {relerr(f32, fo):.2e} against the oracle")
        assert relerr(f64, fo) < 1e-10 and relerr(f32, fo) < 1e-4
        checked += 1 if len(e1g) > 0 else 0
        fresh.close()
        o.close()
    assert ratio < 0.31 and checked >= 3
    g.close()


def test_config4_sweep_512():
    shells = batch_sweep(512, seed=7, loading="pressure")
    meshes = [(s["X"], s["T"]) for s in shells]
    params = [s["params"] for s in shells]
    probe = S.ShellBatch(meshes, params)
    omega = probe.cfl_omega_max(60)
    probe.close()
    params = [dict(p, dt=1.0 / w, dt_max=1.0 / w) for p, w in zip(params, omega)]
    levels = [s["level"] for s in shells]
    say(f"case config 4: {len(shells)} shells, levels {[levels.count(L) for L in (3, 4, 5)]}, pressure ramps")
    b = S.ShellBatch(meshes, params)
    z = np.zeros(3 * int(b.nv.sum()))
    b.set_state(S.ALL_SHELLS, z, v=z, a=z, t=0.0)
    b.apply_initial_conditions()
    pick = [levels.index(L) for L in (3, 4, 5)]
    orcs = {}
    for i in pick:
        X, T = meshes[i]
        o = oracle.Oracle(X, T, **params[i])
        o.set_state(np.zeros(X.size))
        o.init_dynamics()
        orcs[i] = o
    t0 = time.perf_counter()
    for k in range(3):
        reps = b.advance(100)
        for i in pick:
            orcs[i].step(100)
            ug = b.get_state(i)[0]
            uo = orcs[i].get_state()[0]
            say(f"  chunk {k} shell {i} (L{levels[i]}): rel u {relerr(ug, uo):.2e}")
            assert relerr(ug, uo) < 1e-9
    say(f"  300 steps of the batch in {time.perf_counter() - t0:.1f} s (oracle legs included)")
    assert all(r["n_steps"] == 300 and r["capacity_err"] == 0 and r["problem"] < S.STOP_SIM for r in reps)
    U = b.get_state(S.ALL_SHELLS)[0]
    assert np.all(np.isfinite(U))
    b.close()


def test_config5_warm_start_1024():
    from shell_amd.surrogate import Surrogate
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    d = np.load(os.path.join(root, "shell_amd", "data", "warm_L4_4096.npz"))
    conv = d["converged"]
    sg = Surrogate(d["umean"], d["basis"][:64].astype(np.float64), hidden=(256, 256, 256), seed=1)
    for ep, lr in ((100, 1e-3), (50, 3e-4), (50, 1e-4)):
        sg.fit(d["params"][conv], d["coeffs"][:, :64].astype(np.float64), epochs=ep, batch=256, lr=lr)
    P = W.draw_params(1024, 12)
    b, meshes = W.build(P)
    vref = W.reference_volumes(b)
    say("case config 5: warm start of 1024 unseen level-4 shells")
    rw, wall_w, info = W.warm_solve(b, sg, P, vref, 1e-6, 60000, 2000)
    rp, wall_p = W.plain_solve(b, 1e-6, 60000, 2000)
    cw = sum(r["converged_step"] >= 0 for r in rw)
    cp = sum(r["converged_step"] >= 0 for r in rp)
    sw = sum(r["n_steps"] for r in rw)
    sp = sum(r["converged_step"] if r["converged_step"] >= 0 else r["n_steps"] for r in rp)
    say(f"  warm: {cw} converged in {wall_w:.1f} s ({sw} steps), plain: {cp} in {wall_p:.1f} s ({sp} steps), "
        f"fallbacks {len(info['fallback'])}, stopped {sum(S.stopped(r) for r in rw)}")
    assert cw >= 1010 and cp >= 1010
    assert sw < sp and wall_w < wall_p
    b.close()
