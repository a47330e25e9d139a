"""Host side of the benchmark (no GPU): the one-GPU contract of the command
line and the algorithmic byte count of the element kernel, whose block
records and entries are counted here a second way from the oracle's
patches."""
import json
import os
import subprocess
import sys

import numpy as np

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert msg["n_gpus"] == 2 and "one GPU" in msg["error"]


def test_block_records_and_entries_from_the_patches():
    """records = patch nodes of each 256-element block, summed; entries =
    (element, slot) pairs; the patches from the oracle's one-ring lists"""
    X, T = O.icosphere(3)
    o = O.Oracle(X, T)
    idx = o.patches()[0]
    o.close()
    records = sum(len(set(idx[b:b + 256][idx[b:b + 256] >= 0].tolist())) for b in range(0, len(T), 256))
    entries = int((idx >= 0).sum())
    assert bench.force_blocks(T, len(X)) == (records, entries)


def test_element_bytes_per_element_level4():
    """325 B per level-4 element with the displacement trigger, 313 B without"""
    X, T = O.icosphere(4)
    per = bench.element_kernel_bytes(T, len(X), True) / len(T)
    assert abs(per - 324.8) < 0.1
    assert abs(bench.element_kernel_bytes(T, len(X), False) / len(T) - (per - 24 * len(X) / len(T))) < 1e-9
