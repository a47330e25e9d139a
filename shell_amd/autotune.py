"""Launch-shape autotuner of the tunable kernels.

Candidates (shx_shapes): waves per SIMD of the element kernel (3, 4), the
node kernel (4, 6, 8), the contact kernel (2, 3, 4) and the exact-distance
pass (2, 3, 4) -- through their launch bounds, i.e. the registers a thread
may use, traded against spills -- and the pair search's lanes per element
(8, 16, 32).  Results do not depend on the shapes, only the speed.

The tuner times every candidate of one shape at a time on the workload
(coordinate search from the defaults, the best value kept before the next
shape): rest state, apply_initial_conditions (the initial contact search of
every shell) and STEPS step attempts launched one by one with HIP events;
the score of a candidate is the total time of the kernel class it shapes
(launch counts are the same for every candidate, the runs being identical).
The choice is stored per workload in shell_amd/data/shapes.json, which
bench.py reads; the candidate table is printed as markdown.

usage: python -m shell_amd.autotune warm1024|sweep512 [STEPS] [--out PATH]
(--out: the store to update, default shell_amd/data/shapes.json)
"""
import json
import os
import sys
import time

import numpy as np

from . import ALL_SHELLS, DEFAULT_SHAPES, FP64, SHAPE_FIELDS

CANDIDATES = {"elem_waves": (3, 4), "node_waves": (4, 6, 8), "contact_waves": (2, 3, 4),
              "exact_waves": (2, 3, 4), "bp_lanes": (8, 16, 32)}
KERNEL_OF = {"elem_waves": "elem_force", "node_waves": "node_newmark", "contact_waves": "contact_narrow",
             "exact_waves": "bp_exact", "bp_lanes": "bp_pairs"}
STORE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "shapes.json")


def load_shapes(workload, path=STORE):
    """the stored shapes of a workload, or None"""
    try:
        with open(path) as f:
            d = json.load(f)
    except (OSError, ValueError):
        return None
    w = d.get(workload)
    if not w:
        return None
    return {k: int(w["shapes"][k]) for k in SHAPE_FIELDS}


def _run(b, shapes, steps, precision):
    b.set_shapes(shapes)
    z = np.zeros(3 * int(b.nv.sum()))
    b.set_state(ALL_SHELLS, z, v=z, a=z, t=0.0)
    b.set_profiling(True)
    b.apply_initial_conditions(precision)
    b.advance(steps, precision=precision)
    kt = b.kernel_times()
    b.set_profiling(False)
    return kt


def tune(b, steps=100, precision=FP64, log=print):
    """coordinate search over CANDIDATES; returns (shapes, rows), rows =
    (shape, candidate, kernel class, launches, avg ms, total ms, chosen)"""
    best = dict(DEFAULT_SHAPES)
    rows = []
    for dim, cands in CANDIDATES.items():
        cls = KERNEL_OF[dim]
        scores = {}
        for c in cands:
            kt = _run(b, dict(best, **{dim: c}), steps, precision)
            avg, n = kt[cls]
            scores[c] = (avg * n, avg, n)
            log(f"  {dim}={c}: {cls} {n} launches, {avg * 1e3:.1f} us avg, {avg * n:.2f} ms total")
        pick = min(cands, key=lambda c: scores[c][0])
        best[dim] = pick
        for c in cands:
            rows.append((dim, c, cls, scores[c][2], scores[c][1], scores[c][0], c == pick))
    b.set_shapes(best)
    return best, rows


def markdown(rows):
    out = ["| shape | candidate | kernel | launches | avg µs | total ms | chosen |",
           "|---|---:|---|---:|---:|---:|:-:|"]
    for dim, c, cls, n, avg, tot, pick in rows:
        out.append(f"| {dim} | {c} | {cls} | {n} | {avg * 1e3:.1f} | {tot:.2f} | {'yes' if pick else ''} |")
    return "\n".join(out)


def build_workload(name):
    """the batches bench.py measures: the 1024 level-4 warm-start shells
    (parameters seed 12) or the 512-shell sweep (seed 7), dt = 0.5 x CFL"""
    from . import ShellBatch
    if name == "warm1024":
        from . import warmstart as W
        b, _ = W.build(W.draw_params(1024, 12))
        return b
    if name == "sweep512":
        from .workloads import batch_sweep
        shells = batch_sweep(512, seed=7)
        meshes = [(s["X"], s["T"]) for s in shells]
        params = [s["params"] for s in shells]
        probe = ShellBatch(meshes, params)
        omega = probe.cfl_omega_max(60)
        probe.close()
        return ShellBatch(meshes, [dict(p, dt=1.0 / w, dt_max=1.0 / w) for p, w in zip(params, omega)])
    raise ValueError(f"unknown workload {name!r}")


def main(argv=None):
    argv = list(sys.argv[1:] if argv is None else argv)
    store = STORE
    if "--out" in argv:
        k = argv.index("--out")
        store = argv[k + 1]
        del argv[k:k + 2]
    name = argv[0]
    steps = int(argv[1]) if len(argv) > 1 else 100
    t0 = time.perf_counter()
    b = build_workload(name)
    print(f"workload {name}: {b.S} shells, built in {time.perf_counter() - t0:.1f} s", flush=True)
    shapes, rows = tune(b, steps, log=lambda s: print(s, flush=True))
    b.close()
    try:
        with open(store) as f:
            d = json.load(f)
    except (OSError, ValueError):
        d = {}
    d[name] = {"shapes": shapes, "steps": steps,
               "candidates": [dict(zip(("shape", "candidate", "kernel", "launches", "avg_ms", "total_ms", "chosen"),
                                       (r[0], r[1], r[2], int(r[3]), float(r[4]), float(r[5]), bool(r[6]))))
                              for r in rows]}
    with open(store, "w") as f:
        json.dump(d, f, indent=1)
    print(f"chosen {shapes}", flush=True)
    print(markdown(rows), flush=True)


if __name__ == "__main__":
    main()
