"""Mesh preparation and the command-line driver on the CPU (no GPU needed):
the library's C++ icosphere generator writes the reference's mesh files
byte for byte, OFF text round-trips, midpoint refinement and the driver's
node transforms do what the driver's mesh preparation does (shell.C:146-230),
and shellhip_run's argument and config errors end the run before any GPU
work with the reference's messages and exit codes (shell.C:23, 80-144)."""
import hashlib
import os
import subprocess

import numpy as np
import pytest

import shell_amd as S

OFF_SHA = {3: "461df12da867577434b5f4db6d261d4ac66749fa1e3666df9a07814fd7a45033",
           4: "7a0dbb5a8c29168e4c0feac280debc9304cf20065e52390aeb5129e304025622"}


@pytest.mark.parametrize("level", [3, 4])
def test_generator_writes_the_reference_mesh_files(level):
    text = S.write_off(*S.mesh_icosphere(level))
    assert hashlib.sha256(text.encode()).hexdigest() == OFF_SHA[level]


def test_off_round_trip_and_errors():
    X, T = S.mesh_icosphere(2)
    text = S.write_off(X, T)
    X2, T2 = S.read_off("# a comment line\n" + text)
    assert np.array_equal(T2, T) and np.allclose(X2, X, rtol=1e-5, atol=1e-6)
    with pytest.raises(S.ShellError):
        S.read_off("OFF\n4 1 0\n0 0 0\n1 0 0\n0 1 0\n1 1 0\n4 0 1 2 3\n")   # a quad
    with pytest.raises(S.ShellError):
        S.read_off("PLY\n")
    with pytest.raises(S.ShellError):
        S.read_off("OFF\n3 1 0\n0 0 0\n1 0 0\n0 1 0\n3 0 1 7\n")          # vertex id out of range


def test_midpoint_refinement_is_the_next_icosphere_before_projection():
    X3, T3 = S.mesh_icosphere(3, round6=False)
    X4, T4 = S.mesh_icosphere(4, round6=False)
    Xr, Tr = S.refine(X3, T3, 1)
    assert np.array_equal(Tr, T4)                         # same numbering and face order
    assert np.allclose(Xr / np.linalg.norm(Xr, axis=1)[:, None], X4, rtol=0, atol=1e-15)
    # edge midpoints lie on the straight edges
    e = Tr[3]                                             # the inner triangle of the first face
    assert np.allclose(Xr[e[0]], (X3[T3[0, 0]] + X3[T3[0, 1]]) / 2)


def test_node_transforms():
    X, _ = S.mesh_icosphere(2, round6=False)
    P = S.transform(X, project=(2.0, 2.0, 2.0))
    assert np.allclose(np.linalg.norm(P, axis=1), 2.0, rtol=1e-15)
    X = X[np.hypot(X[:, 0], X[:, 1]) > 0.1]              # off the axis of the cylinder
    E = S.transform(X, project=(1.0, 2.0, 0.0))           # elliptic cylinder: x^2 + y^2/4 = 1, z kept
    assert np.allclose(E[:, 0] ** 2 + E[:, 1] ** 2 / 4, 1.0) and np.array_equal(E[:, 2], X[:, 2])
    Sc = S.transform(X, scale=(2.0, 3.0, 0.5), translate=(1.0, -1.0, 0.25))
    assert np.allclose(Sc, X * [2.0, 3.0, 0.5] + [1.0, -1.0, 0.25], rtol=0, atol=1e-15)
    # rotations (Euler z-x-z, in turns): rigid, and a quarter turn of phi maps x to -y
    R = S.transform(X, rotate=(0.1, 0.27, 0.4))
    d0 = np.linalg.norm(X[:, None] - X[None], axis=2)
    d1 = np.linalg.norm(R[:, None] - R[None], axis=2)
    assert np.allclose(d0, d1, rtol=0, atol=1e-14)
    q = S.transform(np.array([[1.0, 0.0, 0.0], [0.0, 0.0, 1.0]]), rotate=(0.25, 0.0, 0.0))
    assert np.allclose(q, [[0.0, -1.0, 0.0], [0.0, 0.0, 1.0]], atol=1e-15)
    q = S.transform(np.array([[0.0, 1.0, 0.0]]), rotate=(0.0, 0.25, 0.0))
    assert np.allclose(q, [[0.0, 0.0, -1.0]], atol=1e-15)
    # whole turns leave the nodes alone (shell.C:217-218 tests fmod(rotate, 1))
    assert np.array_equal(S.transform(X, rotate=(1.0, 2.0, 0.0)), X)


def _run(args, cwd):
    assert os.path.exists(S.DRIVER_PATH), "build the driver first"
    return subprocess.run([S.DRIVER_PATH] + args, cwd=cwd, capture_output=True, text=True, timeout=60)


def test_driver_argument_and_config_errors(tmp_path):
    r = _run(["-h"], tmp_path)
    assert r.returncode == 0 and "Usage:" in r.stdout
    r = _run(["-c", "missing.cfg"], tmp_path)
    assert r.returncode == 3 and "Can't open config file \"missing.cfg\"" in r.stdout
    cfg = tmp_path / "config.cfg"
    cfg.write_text("shell_young_modulus2 = 1e9\nshell_shear_modulus = 3.5e8\nextra_order = 1\n")
    r = _run([], tmp_path)
    assert r.returncode == 3 and "quadrature" in r.stdout
    r = _run(["-extra_order", "0"], tmp_path)            # no mesh given
    assert r.returncode == 3 and "-m meshfile.off" in r.stdout
    r = _run(["-extra_order", "0", "-m", "nope.off"], tmp_path)
    assert r.returncode == 3 and "Can't open mesh file" in r.stdout
    # the config is copied into the data directory, as shell.C:105 does
    assert (tmp_path / "data" / "config.cfg").exists()
