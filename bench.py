#!/usr/bin/env python3
"""Benchmark: post-buckled equilibria per second of the explicit
force-and-step path on one MI355X, with the element-force kernel's HBM
roofline and the CPU oracle timed on the host beside it.

A benchmark "step" is one complete equilibrium solve of the workload: a
level-4 icosphere shell (2562 nodes, 5120 Loop elements) deflated by a
prescribed enclosed-volume reduction to 90% (deep in the post-buckled
range: the sphere buckles after about 1.3% volume loss), relaxed by the
damped explicit method until the residual of the deformation
||f - M a_rigid|| / ||f_ext|| <= 1e-6.  Inputs (mesh, tables, state)
are resident in HBM before the timed region; each solve restarts from the
same initial state.  value = converged solves / wall seconds.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import shell_amd as S  # noqa: E402
from shell_amd.workloads import icosphere, p_critical, radial_imperfection  # noqa: E402

HBM_PEAK_GBS = 8000.0


def workload(level):
    """the single-GPU workload (level-4 single shell) and its parameters"""
    h, E1, E2, nu, G12 = 0.01, 1.0, 0.5, 0.3, 0.35
    X, T = icosphere(level)
    X = radial_imperfection(X, 0.01 * h, seed=2024)
    pcr = p_critical(E1, h, nu)
    params = dict(E1=E1, E2=E2, nu=nu, G12=G12, thickness=h, rho=1.0, damping_alpha=0.05,
                  contact=1, mu_s=0.3, mu_d=0.3, cell_size_mult=4.0)
    return X, T, params, pcr


def element_kernel_bytes(ne, nv):
    """algorithmic HBM bytes of one element-force launch (FP64):
    per element: 12 int32 patch ids + 1 valence byte + 26 FP64 table rows
    (R 2x2, JxW, A1 A2 A11 A22 A12, ff1, ff2) read; 12 slot force
    contributions (36 FP64) and JxW a3 (3 FP64) written; element area read
    and written;
    per node: u, v, a (72 B) and the mesh position (24 B) read once."""
    per_elem = 12 * 4 + 1 + 26 * 8 + 39 * 8 + 2 * 8
    per_node = 4 * 3 * 8
    return ne * per_elem + nv * per_node


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2, help="timed equilibrium solves")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--level", type=int, default=4)
    ap.add_argument("--max-steps", type=int, default=400000)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    X, T, prm, pcr = workload(args.level)
    # stable step: half the central-difference limit 2/omega_max of M^-1 K
    probe = S.ShellSystem(X, T, **prm)
    omega = probe.cfl_omega_max(200)
    probe.close()
    dt = 0.5 * 2.0 / omega
    ramp_steps = 5000
    run = dict(prm, bulk_modulus=0.05, volume_final_ratio=0.9, volume_ramp_time=ramp_steps * dt,
               volume_at_prediction=1, dt=dt, dt_max=dt)
    g = S.ShellSystem(X, T, **run)
    z = np.zeros(3 * len(X))

    def solve():
        g.set_state(z, z, z, 0.0)
        g.apply_initial_conditions()
        return g.advance(args.max_steps, tol=args.tol, stop_on_converged=True)

    for _ in range(args.warmup):
        solve()
    reps = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        reps.append(solve())
    t1 = time.perf_counter()
    wall = t1 - t0
    converged = sum(1 for r in reps if r["converged_step"] >= 0)
    steps_per_solve = float(np.mean([r["n_steps"] for r in reps]))
    attempts_per_solve = steps_per_solve + float(np.mean([r["n_rejected"] for r in reps]))

    # roofline of the element-force kernel: HIP events on the library's stream
    g.set_state(z, z, z, 0.0)
    g.apply_initial_conditions()
    g.set_profiling(True)
    g.advance(2000)
    kt = g.kernel_times()
    g.set_profiling(False)
    elem_ms = kt["elem_force"][0]
    nbytes = element_kernel_bytes(len(T), len(X))
    achieved = nbytes / (elem_ms * 1e-3) / 1e9 if elem_ms > 0 else 0.0

    cpu = None
    if not args.no_cpu:
        import oracle  # CPU baseline leg only
        threads = os.cpu_count() or 1
        o = oracle.Oracle(X, T, **dict(run, omp_faithful=1))
        o.set_state(z)
        o.init_dynamics()
        n, c0 = 0, time.perf_counter()
        while time.perf_counter() - c0 < args.cpu_seconds:
            o.step(100)
            n += 100
        cpu_wall = time.perf_counter() - c0
        cpu_steps_s = n / cpu_wall
        cpu = {"value": cpu_steps_s / attempts_per_solve, "unit": "equilibria/s", "cores": threads,
               "kind": "port",
               "sample": f"{n} steps of the same workload from the same initial state on the CPU oracle "
                         f"(OpenMP element loops, critical-section assembly), {cpu_wall:.1f} s; "
                         f"equilibria/s = steps/s / {attempts_per_solve:.0f} step attempts per equilibrium "
                         f"measured on the GPU",
               "steps_per_s": cpu_steps_s}

    ms_per_solve = wall / max(args.steps, 1) * 1e3
    out = {
        "metric": "post-buckled equilibria/s at rel residual 1e-6 (plain FP64 GPU method); force GB/s",
        "value": converged / wall if wall > 0 else 0.0,
        "unit": "equilibria/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_solve,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: level-4 icosphere, seeded radial imperfection 0.01h",
        "config": {"workload": f"level-{args.level} icosphere shell, h=0.01, E1=1, E2=0.5, nu12=0.3, G12=0.35, "
                               f"imperfection 0.01h, deflation to V/V0=0.9 over {ramp_steps} steps (bulk modulus "
                               f"0.05), self-contact mu=0.3, dt=0.5*CFL={dt:.4g}, mass damping 0.05, relaxation to "
                               f"deformation residual {args.tol:g}",
                   "nodes": len(X), "elements": len(T), "converged": converged,
                   "solver_steps_per_equilibrium": steps_per_solve,
                   "step_attempts_per_equilibrium": attempts_per_solve,
                   "solver_steps_per_s": sum(r["n_steps"] + r["n_rejected"] for r in reps) / wall},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_elem (element force)", "bytes_per_launch": nbytes,
                     "avg_launch_ms": elem_ms, "kernel_ms": {k: v[0] for k, v in kt.items()}},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
