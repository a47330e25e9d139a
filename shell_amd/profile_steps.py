"""Run a fixed number of explicit step attempts of the benchmark workloads,
for profilers (rocprofv3 --kernel-trace / --pmc): the level-4 single shell,
the seeded batched sweep, or the 1024 level-4 shells of the warm-start test
set (plain start, dt = 0.5 x their CFL bound), FP64 or FP32 split, attempts
launched one by one (no graph); `smooth`: a level-4 shell pancaked to
3 % of its height whose sheets close at 0.04 per unit time, with smooth
(limit-surface) contact and friction, adaptive dt; and `open`: 64 open
square plates of 64 x 64 squares (8192 elements each, with their ghost
layers) pinned on their edges and sagging under their own weight.

usage: python -m shell_amd.profile_steps single|batch|surrogate|smooth|open STEPS [fp32]
"""
import sys

import numpy as np

import shell_amd as S
from shell_amd.workloads import batch_sweep, icosphere, radial_imperfection

# stable steps of the sweep's meshes (half the power-iteration CFL bound of
# the unit icosphere with E1 = 1, rho = 1: omega_max 8.6 / 17.9 / 37.0)
DT = {3: 0.5 * 2.0 / 8.6, 4: 0.5 * 2.0 / 17.9, 5: 0.5 * 2.0 / 37.0}


def main():
    what, steps = sys.argv[1], int(sys.argv[2])
    prec = S.FP32_SPLIT if len(sys.argv) > 3 and sys.argv[3] == "fp32" else S.FP64
    if what == "single":
        X, T = icosphere(4)
        X = radial_imperfection(X, 1e-4, seed=2024)
        p = dict(thickness=0.01, E2=0.5, damping_alpha=0.05, contact=1, mu_s=0.3, mu_d=0.3, cell_size_mult=4.0,
                 cell_margin=0.1, rebuild_trigger=1, bulk_modulus=0.05, volume_final_ratio=0.9,
                 volume_ramp_time=5000 * DT[4],
                 volume_at_prediction=1, dt=DT[4], dt_max=DT[4])
        meshes, params = [(X, T)], [p]
    elif what == "surrogate":
        from shell_amd import warmstart as W
        b, meshes = W.build(W.draw_params(1024, 12))
        params = None
    elif what == "smooth":
        X, T = icosphere(4)
        X = radial_imperfection(X, 1e-4, seed=7)
        Y = X.copy()
        Y[:, 2] *= 0.03
        u0 = (Y - X).reshape(-1)
        v0 = np.zeros_like(X)
        v0[:, 2] = -0.02 * np.sign(X[:, 2])
        v0 = v0.reshape(-1)
        meshes = [(X, T)]
        params = [dict(contact=2, mu_s=0.3, mu_d=0.3, cell_size_mult=1.0, dt=0.01, dt_max=0.01, damping_alpha=0.05,
                       adapt_dt_every=1)]
    elif what == "open":
        n = 64
        i, j = np.meshgrid(np.arange(n + 1), np.arange(n + 1), indexing="ij")
        X = np.stack([i.ravel() / n, np.zeros(i.size), j.ravel() / n], axis=1)     # the x-z plane
        q = (i[:-1, :-1] * (n + 1) + j[:-1, :-1]).ravel()
        T = np.concatenate([np.stack([q, q + n + 1, q + n + 2], 1), np.stack([q, q + n + 2, q + 1], 1)])
        Xg, Tg, ng = S.mesh_add_ghosts(X, T.astype(np.int32))
        meshes = [(Xg, Tg)] * 64
        params = [dict(E2=1.0, G12=1.0 / 2.6, thickness=0.05, gravity=3e-3, gravity_dir=1, damping_alpha=0.1,
                       ghost_elems=ng, shell_boundary=2, bc_every=1, dt=0.0018)] * 64
    else:
        shells = batch_sweep(512, seed=7)
        meshes = [(s["X"], s["T"]) for s in shells]
        params = [dict(s["params"], dt=DT[s["level"]], dt_max=DT[s["level"]]) for s in shells]
    if params is not None:
        b = S.ShellBatch(meshes, params)
    b.set_graphs(False)
    z = np.zeros(3 * int(b.nv.sum()))
    if what == "smooth":
        b.set_state(0, u0, v=v0, a=z, t=0.0)
    else:
        b.set_state(S.ALL_SHELLS, z, v=z, a=z, t=0.0)
    b.apply_initial_conditions(precision=prec)
    reps = b.advance(steps, precision=prec)
    print(f"{what}: {steps} steps, {sum(r['n_rebuilds'] for r in reps)} rebuilds, "
          f"{sum(len(T) for _, T in meshes)} elements", flush=True)


if __name__ == "__main__":
    main()
