"""Learned warm start of the batched damped relaxation, level-4 shells.

Workload family: level-4 icosphere shells (2562 nodes, 5120 triangles), one
fixed radial imperfection pattern (unit Gaussian per vertex, seed
PATTERN_SEED; an inward Gaussian dimple is the alternative pattern) scaled
per shell, deflation by the volume constraint to V/V0 at
the end of a RAMP_TIME ramp, self-contact with Coulomb friction (contact
lists rebuilt on the displacement trigger), mass damping 0.05, dt = 0.5 x
the shell's CFL bound.  Six parameters per shell (PARAMS):

    E2/E1 ~ U[0.3, 1], nu12 ~ U[0.25, 0.35], h/R ~ logU[0.008, 0.016],
    V/V0 ~ U[0.86, 0.92], mu ~ U[0, 0.5], imperfection amplitude / h ~ U[0.005, 0.02]

(the load parameter is the prescribed volume ratio: the shells are loaded
by deflation, under which post-buckled equilibria are stable).

Plain start: u = 0 at t = 0, the load ramps up, then the relaxation runs
until the deformation residual ||f - M a_rigid|| / ||f_ext|| <= tol.

Warm start: the surrogate's displacement at t = RAMP_TIME (full load),
v = 0, the volume constraint's V0 fixed to the undeformed shell's volume;
QUENCH_STEPS steps with extra mass damping QUENCH / dt remove the
prediction's high-frequency error (a small displacement error in the stiff
membrane modes gives a residual of order 10), then the same relaxation as
the plain start judges the state.  A shell falls back to the plain start
when its warm state interpenetrates (CONTACT_PROBLEM at the initial contact
search), overflows a contact list, or when its residual after the quench
and PROBE_STEPS more steps is not below the residual of the warm state.
"""
import sys
import time

import numpy as np

from . import ALL_SHELLS, E_CAPACITY, FP64, ShellBatch, ShellError, stopped
from .workloads import icosphere

LEVEL = 4
PATTERN = "noise"
PATTERN_SEED = 4242
DIMPLE_WIDTH = 0.15
RAMP_TIME = 280.0
PARAMS = ("E2/E1", "nu12", "h/R", "V/V0", "mu", "imperfection/h")
CONTACT_PROBLEM = 6
PROBE_STEPS = 1000
QUENCH_STEPS = 2000
QUENCH = 0.1


def log(*a):
    print(*a, file=sys.stderr, flush=True)


AMPLITUDE = {"dimple": (0.1, 0.5), "noise": (0.005, 0.02)}


def draw_params(n, seed, kind=None):
    """n x 6 parameter sets (PARAMS order) from a seed"""
    rng = np.random.default_rng(seed)
    P = np.empty((n, 6))
    P[:, 0] = rng.uniform(0.3, 1.0, n)
    P[:, 1] = rng.uniform(0.25, 0.35, n)
    P[:, 2] = np.exp(rng.uniform(np.log(0.008), np.log(0.016), n))
    P[:, 3] = rng.uniform(0.86, 0.92, n)
    P[:, 4] = rng.uniform(0.0, 0.5, n)
    P[:, 5] = rng.uniform(*AMPLITUDE[kind or PATTERN], n)
    return P


def pattern(X0, kind=None):
    """the fixed imperfection pattern (radial amplitude per vertex, unit
    scale): "dimple", an inward Gaussian dent of angular width DIMPLE_WIDTH
    centred on the equator at +x; "noise", unit Gaussian noise per vertex"""
    kind = kind or PATTERN
    n = X0 / np.linalg.norm(X0, axis=1)[:, None]
    if kind == "noise":
        return np.random.default_rng(PATTERN_SEED).normal(size=(len(X0), 1))
    th = np.arccos(np.clip(n[:, 0], -1.0, 1.0))
    return -np.exp(-(th / DIMPLE_WIDTH) ** 2)[:, None]


def shells(P, kind=None):
    """meshes and parameter dicts of the shells of parameter rows P"""
    X0, T = icosphere(LEVEL)
    n = X0 / np.linalg.norm(X0, axis=1)[:, None]
    g = pattern(X0, kind)
    meshes, prms = [], []
    for e2, nu, h, ratio, mu, amp in np.asarray(P, np.float64):
        meshes.append((X0 + n * g * (amp * h), T))
        prms.append(dict(E1=1.0, E2=float(e2), nu=float(nu), G12=0.35, thickness=float(h), rho=1.0,
                         damping_alpha=0.05, bulk_modulus=0.05, volume_final_ratio=float(ratio),
                         volume_ramp_time=RAMP_TIME, volume_at_prediction=1, contact=1, mu_s=float(mu),
                         mu_d=float(mu), cell_size_mult=4.0, cell_margin=0.1, rebuild_trigger=1))
    return meshes, prms


def build(P, cfl_iters=60, kind=None):
    """a batch of the shells of P with dt = 0.5 * 2 / omega_max per shell"""
    meshes, prms = shells(P, kind)
    probe = ShellBatch(meshes, prms)
    omega = probe.cfl_omega_max(cfl_iters)
    probe.close()
    prms = [dict(p, dt=1.0 / w, dt_max=1.0 / w) for p, w in zip(prms, omega)]
    return ShellBatch(meshes, prms), meshes


def _converged_or_dead(reps):
    return all(r["converged_step"] >= 0 or stopped(r) for r in reps)


def _advance(b, n, tol, precision, tolerant):
    """n steps (stop on convergence); tolerant: a shell whose device lists
    overflow stops with capacity_err set and the others go on"""
    try:
        return b.advance(n, tol=tol, stop_on_converged=True, precision=precision)
    except ShellError as e:
        if not tolerant or e.code != E_CAPACITY:
            raise
        return b.reports()


def _relax(b, tol, max_steps, chunk, t0, label, precision=FP64, on_overflow=None):
    """step the batch until every shell has converged (or max_steps);
    on_overflow(shells) restarts shells whose lists overflowed"""
    done, reps = 0, None
    while done < max_steps:
        n = min(chunk, max_steps - done)
        reps = _advance(b, n, tol, precision, on_overflow is not None)
        over = [i for i, r in enumerate(reps) if r["capacity_err"]]
        if over:
            on_overflow(over)
            reps = b.reports()
        done += n
        nconv = sum(r["converged_step"] >= 0 for r in reps)
        log(f"  {label} {done} steps {time.perf_counter() - t0:.1f} s converged {nconv}/{len(reps)}")
        if _converged_or_dead(reps):
            break
    return reps


def _rest(b):
    """u = v = a = 0 at t = 0 for every shell"""
    z = np.zeros(3 * int(b.nv.sum()))
    b.set_state(ALL_SHELLS, z, v=z, a=z, t=0.0)


def plain_solve(b, tol=1e-6, max_steps=60000, chunk=2000, precision=FP64):
    """plain start of every shell; returns (reports, wall seconds)"""
    t0 = time.perf_counter()
    _rest(b)
    b.apply_initial_conditions(precision)
    reps = _relax(b, tol, max_steps, chunk, t0, "plain", precision)
    return reps, time.perf_counter() - t0


def reference_volumes(b):
    """enclosed volume of every shell's undeformed (u = 0) state"""
    _rest(b)
    b.apply_initial_conditions()
    return np.array([r["volume"] for r in b.reports()])


def warm_solve(b, surrogate, P, vref, tol=1e-6, max_steps=60000, chunk=2000, probe_steps=PROBE_STEPS,
               precision=FP64, quench_steps=QUENCH_STEPS, quench=QUENCH):
    """warm start of every shell from the surrogate's prediction, with the
    fallback to a plain start; returns (reports, wall seconds, info).  vref:
    the shells' undeformed volumes (reference_volumes, computed before the
    timed region: they depend on the mesh only).  precision: SHX_FP64 or
    SHX_FP32_SPLIT (FP32 element forces until the residual is within 1000 tol,
    then FP64).  quench_steps > 0: the warm-started shells first take that
    many steps with extra mass damping quench / dt (about critical for the
    highest mode: a fraction `quench` of 1/dt), which removes most of the
    prediction's high-frequency error; a shell fails the probe when its
    residual after the quench and probe_steps more steps is not below the
    residual of its warm state."""
    t0 = time.perf_counter()
    S = b.S
    U0 = surrogate.predict(P)
    t_pred = time.perf_counter() - t0
    b.set_state(ALL_SHELLS, U0.reshape(-1), v=np.zeros(U0.size), a=np.zeros(U0.size), t=RAMP_TIME)
    try:
        b.init_shells(np.arange(S), volume0=vref, precision=precision)
    except ShellError as e:
        if e.code != E_CAPACITY:
            raise
    reps = b.reports()
    r0 = np.array([r["residual_deform"] for r in reps])
    bad = [i for i, r in enumerate(reps)
           if r["problem"] == CONTACT_PROBLEM or r["capacity_err"] or not np.isfinite(r0[i])]
    fell = set(bad)
    late = []

    def fallback(idx):
        if not idx:
            return
        for i in idx:
            z = np.zeros(3 * int(b.nv[i]))
            b.set_state(i, z, v=z, a=z, t=0.0)
        b.init_shells(idx, precision=precision)

    def overflow(idx):
        # lists that overflow during the relaxation: the warm state was not
        # usable; a plain start that overflows is a real capacity error
        plain = [i for i in idx if i in fell]
        if plain:
            raise ShellError(f"device list capacity exceeded in plain-started shells {plain}", E_CAPACITY)
        fell.update(idx)
        late.extend(idx)
        fallback(idx)

    fallback(bad)
    rq = r0
    if quench_steps > 0:
        dt = np.array([b.params[i].dt for i in range(S)])
        alpha = np.where([i in fell for i in range(S)], 0.0, quench / dt)
        b.set_extra_damping(alpha)
        reps = _advance(b, quench_steps, tol, precision, True)
        b.set_extra_damping(None)
        rq = np.array([r["residual_deform"] for r in reps])
        over = [i for i, r in enumerate(reps) if r["capacity_err"] and i not in fell]
        fallback(over)
        fell.update(over)
    reps = _advance(b, probe_steps, tol, precision, True)
    grew = [i for i, r in enumerate(reps)
            if i not in fell and (r["capacity_err"] or (r["converged_step"] < 0 and not r["residual_deform"] < r0[i]))]
    fallback(grew)
    fell.update(grew)
    log(f"  warm: {len(bad)} interpenetrating, {len(grew)} with a growing residual after {probe_steps} steps")
    reps = b.reports()
    if not _converged_or_dead(reps):
        reps = _relax(b, tol, max_steps, chunk, t0, "warm", precision, on_overflow=overflow)
    wall = time.perf_counter() - t0
    info = {"steps_to_converge": [r["converged_step"] for r in reps], "fallback_interpenetrating": len(bad), "fallback_residual_growth": len(grew),
            "fallback_overflow_later": len(late),
            "fallback": sorted(fell), "predict_s": t_pred, "r0_median": float(np.median(r0[np.isfinite(r0)])),
            "r_quench_median": float(np.median(rq[np.isfinite(rq)]))}
    return reps, wall, info


def stream_solve(b, P, vref, slots, surrogate=None, tol=1e-6, max_steps=60000, chunk=500,
                 probe_steps=PROBE_STEPS, precision=FP64):
    """Streaming batch: the batch's shells form a queue of which at most
    `slots` step at once; when a shell converges, stops or reaches max_steps,
    the next shell of the queue starts in its place.  surrogate given: a new
    shell starts from the prediction (t = RAMP_TIME, V0 = vref) and falls
    back to a plain start when its warm state interpenetrates or overflows a
    list, or when its residual after probe_steps steps is not below its
    starting residual; otherwise shells start from rest.  No quench (the
    relaxation is expected to carry the kinetic damping).
    Returns (reports, wall seconds, info); reports count steps since each
    shell's (last) start."""
    t0 = time.perf_counter()
    S = b.S
    queue = list(range(S))
    active, warm, r0 = set(), {}, {}
    fell = []
    b.park(np.arange(S))

    def start(idx, use_warm):
        if not idx:
            return
        if use_warm:
            U0 = surrogate.predict(P[idx])
            for k, i in enumerate(idx):
                z = np.zeros(U0.shape[1])
                b.set_state(i, U0[k], v=z, a=z, t=RAMP_TIME)
        else:
            for i in idx:
                z = np.zeros(3 * int(b.nv[i]))
                b.set_state(i, z, v=z, a=z, t=0.0)
        try:
            b.init_shells(idx, volume0=vref[idx] if use_warm else None, precision=precision)
        except ShellError as e:
            if e.code != E_CAPACITY:
                raise
        reps = b.reports()
        for i in idx:
            warm[i] = use_warm
            r0[i] = reps[i]["residual_deform"]
            active.add(i)
        if use_warm:
            bad = [i for i in idx if reps[i]["problem"] == CONTACT_PROBLEM or reps[i]["capacity_err"]
                   or not np.isfinite(r0[i])]
            fell.extend(bad)
            start(bad, False)

    def refill():
        n = min(slots - len(active), len(queue))
        idx = [queue.pop(0) for _ in range(n)]
        start(idx, surrogate is not None)

    refill()
    final = [None] * S
    while active:
        reps = _advance(b, chunk, tol, precision, True)
        finished, back = [], []
        for i in sorted(active):
            r = reps[i]
            if r["converged_step"] >= 0 or stopped(r) or r["n_steps"] >= max_steps:
                finished.append(i)
            elif r["capacity_err"]:
                if warm[i]:
                    back.append(i)
                else:
                    finished.append(i)
            elif warm[i] and r["n_steps"] >= probe_steps and r["n_steps"] - chunk < probe_steps \
                    and not r["residual_deform"] < r0[i]:
                back.append(i)
        for i in finished:
            active.discard(i)
            final[i] = r_copy = dict(reps[i])
            r_copy["warm"] = warm[i]
        if finished:
            b.park(finished)
        for i in back:
            active.discard(i)
        fell.extend(back)
        start(back, False)
        refill()
    wall = time.perf_counter() - t0
    return final, wall, {"fallback": sorted(set(fell))}


def align(Xs, U):
    """remove the rigid motion of each equilibrium: u' such that X + u' is
    X + u translated and rotated (Kabsch) onto X; Xs: one nv x 3 reference
    per sample, U: samples x 3nv"""
    out = np.empty_like(U)
    for k, u in enumerate(U):
        if not np.all(np.isfinite(u)):
            out[k] = np.nan
            continue
        X = Xs[k]
        c0 = X.mean(0)
        x = X + u.reshape(-1, 3)
        c = x.mean(0)
        H = (x - c).T @ (X - c0)
        W, _, Vt = np.linalg.svd(H)
        d = np.sign(np.linalg.det(W @ Vt))
        R = (W @ np.diag([1.0, 1.0, d]) @ Vt).T          # R (x - c) ~ X - c0
        out[k] = ((x - c) @ R.T + c0 - X).reshape(-1)
    return out


def generate(n, seed, batch=1024, tol=1e-6, max_steps=60000, chunk=2000, k=64, kind=None, save_u=False):
    """solve n plain-started shells (in batches), align their equilibria and
    reduce them to k principal components; returns the training-set dict"""
    from .surrogate import pca
    kind = kind or PATTERN
    P = draw_params(n, seed, kind)
    U, steps, conv, walls = [], [], [], []
    for s in range(0, n, batch):
        Pb = P[s:s + batch]
        t0 = time.perf_counter()
        b, meshes = build(Pb, kind=kind)
        log(f"batch {s // batch}: {len(Pb)} shells built in {time.perf_counter() - t0:.1f} s")
        reps, wall = plain_solve(b, tol, max_steps, chunk)
        walls.append(wall)
        u, _, _, _ = b.get_state(ALL_SHELLS)
        u = u.reshape(len(Pb), -1)
        U.append(align([X for X, _ in meshes], u))
        steps += [r["converged_step"] for r in reps]
        conv += [r["converged_step"] >= 0 for r in reps]
        b.close()
    U = np.concatenate(U)
    conv = np.array(conv)
    umean, basis, C, explained = pca(U[conv], k)
    nrm = np.linalg.norm(U[conv] - umean)
    rel = {f"pca_rel_err_{q}": float(np.linalg.norm(U[conv] - umean - C[:, :q] @ basis[:q]) / nrm)
           for q in (8, 16, 32, 64, 128) if q <= len(basis)}
    resid = U[conv] - umean - C @ basis
    extra = {"U": U.astype(np.float32)} if save_u else {}
    return {**rel, **extra, "params": P, "converged": conv, "steps": np.array(steps), "umean": umean,
            "basis": basis.astype(np.float32), "coeffs": C.astype(np.float32), "explained": explained,
            "pca_rel_err": float(np.linalg.norm(resid) / np.linalg.norm(U[conv])), "solve_s": np.array(walls),
            "tol": tol, "max_steps": max_steps, "seed": seed, "pattern": kind}


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(description="generate the warm-start training set (plain solves on the GPU)")
    ap.add_argument("--n", type=int, default=4096)
    ap.add_argument("--seed", type=int, default=11)
    ap.add_argument("--batch", type=int, default=1024)
    ap.add_argument("--max-steps", type=int, default=60000)
    ap.add_argument("--k", type=int, default=128)
    ap.add_argument("--pattern", choices=["dimple", "noise"], default=PATTERN)
    ap.add_argument("--save-u", action="store_true", help="also store the aligned equilibria (float32)")
    ap.add_argument("--out", required=True)
    a = ap.parse_args(argv)
    d = generate(a.n, a.seed, a.batch, max_steps=a.max_steps, k=a.k, kind=a.pattern, save_u=a.save_u)
    np.savez_compressed(a.out, **d)
    log(f"{int(d['converged'].sum())}/{a.n} converged; PCA keeps {d['explained']:.6f} of the variance, "
        f"relative reconstruction error {d['pca_rel_err']:.3e}; solve {d['solve_s'].sum():.1f} s")
    log("  " + ", ".join(f"{k} {v:.3e}" for k, v in d.items() if k.startswith("pca_rel_err_")))


if __name__ == "__main__":
    main()
