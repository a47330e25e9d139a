"""Writes the golden fixtures of tests/test_golden.py with the CPU FP64
oracle (the reference itself cannot be built or run here: libMesh and
muparser are absent and Contact.C is missing).  Run from the repository
root:  python tests/golden/make_golden.py

golden_L3.npz
  basis_v{3,4,5,6}  Loop basis at the quadrature point (6 x (N+6))
  X, T              level-3 icosphere (the reference's icosasphere3.off) with
                    a seeded radial imperfection of 0.01 h (seed 3)
  u, f, fext, en    forces and energies (E_tens, E_bend) at a seeded random
                    displacement (std 1e-3), external pressure -2 p_cr, t = 0.5
  pe1, pe2          sorted contact candidates of the shell squashed to 0.6%
                    of its height (mu 0.3, cell_size_mult 1)
  traj_u            displacement after 100 constant-dt steps from u0 = 1e-4 N(0,1)
  dyn_u, dyn_rep    squashed shell closing at 0.02 with contact, friction and
                    adaptive dt: displacement, rebuilds and rejections after
                    100 steps
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import oracle  # noqa: E402
from conftest import imperfect  # noqa: E402

H = 0.01
PCR = 2 * H * H / np.sqrt(3 * (1 - 0.3 ** 2))

FORCE_KW = dict(pressure_final=-2 * PCR)
PAIR_KW = dict(contact=1, mu_s=0.3, mu_d=0.3, cell_size_mult=1.0)
TRAJ_KW = dict(pressure_final=-2 * PCR, pressure_ramp_time=5.0, dt=0.1, damping_alpha=0.05)
DYN_KW = dict(dt=0.02, dt_max=0.02, damping_alpha=0.05, adapt_dt_every=1, contact=1, mu_s=0.3, mu_d=0.3,
              cell_size_mult=1.0)


def squash(X, s):
    Y = X.copy()
    Y[:, 2] *= s
    return (Y - X).reshape(-1)


def make():
    out = {}
    for v in (3, 4, 5, 6):
        out[f"basis_v{v}"] = oracle.raw_elem_basis(v)
    X0, T = oracle.icosphere(3)
    X = imperfect(X0, 0.01 * H, seed=3)
    out["X"], out["T"] = X, T
    o = oracle.Oracle(X, T, **FORCE_KW)
    u = np.random.default_rng(17).normal(size=X.size) * 1e-3
    f, fe, en, _ = o.forces(u, t=0.5)
    out["u"], out["f"], out["fext"], out["en"] = u, f, fe, en[:2]
    o.close()
    o = oracle.Oracle(X0, T, **PAIR_KW)
    us = squash(X0, 0.006)
    o.set_state(us)
    o.rebuild(force=True)
    out["pe1"], out["pe2"] = o.pairs()
    o.close()
    o = oracle.Oracle(X, T, **TRAJ_KW)
    u0 = np.random.default_rng(5).normal(size=X.size) * 1e-4
    o.set_state(u0)
    o.init_dynamics()
    o.step(100)
    out["traj_u0"], out["traj_u"] = u0, o.get_state()[0]
    o.close()
    o = oracle.Oracle(X, T, **DYN_KW)
    ud = squash(X, 0.03)
    vd = np.zeros_like(ud).reshape(-1, 3)
    vd[:, 2] = -0.02 * np.sign(X[:, 2])
    o.set_state(ud, vd.reshape(-1))
    o.init_dynamics()
    r = o.step(100)
    out["dyn_u0"], out["dyn_v0"], out["dyn_u"] = ud, vd.reshape(-1), o.get_state()[0]
    out["dyn_rep"] = np.array([r["n_rebuilds"], r["n_rejected"], r["n_pairs"]], dtype=np.int64)
    o.close()
    return out


if __name__ == "__main__":
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden_L3.npz")
    np.savez_compressed(path, **make())
    print("wrote", path, os.path.getsize(path), "bytes")
