"""The C-ABI boundary on the CPU (no GPU needed): libshellhip.so loads and
exports every entry point include/shellhip.h declares, the ctypes mirror of
its structs matches the C layout (checked with a probe compiled by gcc from
the header), and the host-compiled copies of the device's narrow-phase and
basis routines agree with the oracle."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import oracle
import shell_amd as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "shellhip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(shx_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_entry_point():
    names = declared_functions()
    assert len(names) >= 25
    L = S.lib()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(S.EXPORTS) <= set(names)


def test_struct_layout_matches_header(tmp_path):
    src = tmp_path / "probe.c"
    fields_p = [f for f, _ in S.Params._fields_]
    fields_r = [f for f, _ in S.Report._fields_]
    fields_c = [f for f, _ in S.Config._fields_]
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "shellhip.h"', 'int main(void) {',
             'printf("%zu %zu %zu\\n", sizeof(shx_params), sizeof(shx_report), sizeof(shx_config));']
    lines += [f'printf("p {f} %zu\\n", offsetof(shx_params, {f}));' for f in fields_p]
    lines += [f'printf("r {f} %zu\\n", offsetof(shx_report, {f}));' for f in fields_r]
    lines += [f'printf("c {f} %zu\\n", offsetof(shx_config, {f}));' for f in fields_c]
    lines += ['return 0; }']
    src.write_text("\n".join(lines))
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    sp, sr, sc = (int(x) for x in out[0].split())
    assert sp == ctypes.sizeof(S.Params) and sr == ctypes.sizeof(S.Report) and sc == ctypes.sizeof(S.Config)
    off = {}
    for ln in out[1:]:
        if ln:
            k, f, o = ln.split()
            off[(k, f)] = int(o)
    for f in fields_p:
        assert off[("p", f)] == getattr(S.Params, f).offset, f
    for f in fields_r:
        assert off[("r", f)] == getattr(S.Report, f).offset, f
    for f in fields_c:
        assert off[("c", f)] == getattr(S.Config, f).offset, f


def test_host_triangle_distance_bitexact_with_oracle():
    """the library's narrow phase (the same source the GPU runs, host-compiled
    with -ffp-contract=off) rounds like the oracle's on random and on
    degenerate (parallel, coplanar, touching) triangle pairs"""
    rng = np.random.default_rng(11)
    cases = [rng.normal(size=18) for _ in range(4000)]
    base = rng.normal(size=9)
    cases += [np.concatenate([base, base + np.array([0, 0, 0.3] * 3)])]          # parallel copy
    cases += [np.concatenate([base, base[[3, 4, 5, 6, 7, 8, 0, 1, 2]] + 1e-3])]  # nearly coincident
    flat = rng.normal(size=18)
    flat[2::3] = 0.0                                                              # coplanar pair
    cases.append(flat)
    for P in cases:
        g = S.tri_distance(P)
        o = oracle.tri_distance(P)
        assert np.array_equal(g, o), (P, g, o)


@pytest.mark.parametrize("valence", [3, 4, 5, 6])
def test_host_element_basis_matches_oracle(valence):
    """two independent constructions of the Loop basis at the quadrature
    point (monomial expansion in the library, Stam box-spline terms with the
    subdivision rows in the oracle)"""
    g = S.raw_elem_basis(valence)
    o = oracle.raw_elem_basis(valence)
    assert g.shape == o.shape
    assert np.max(np.abs(g - o)) < 1e-13


def test_library_refuses_bad_input_without_gpu():
    """argument checks happen before any device call"""
    h = ctypes.c_void_p()
    X = np.zeros((3, 3))
    T = np.array([[0, 1, 2]], dtype=np.int32)
    p = S.make_params(thickness=-1.0)
    r = S.lib().shx_create(ctypes.byref(h), 3, 1, S._dp(X), T.ctypes.data_as(S._I), ctypes.byref(p), 0)
    assert r == -1
    assert b"invalid parameters" in S.lib().shx_last_error(None)
