"""The command-line driver shellhip_run (a C++ program over the C-ABI, the
reference's shell.C) end to end on the GPU: config file and overrides, OFF
mesh, initial distortion, output cadence (property lines every data_time,
.vtu files every vtk_time, state files), a restart from a state file that
continues the run exactly, and the same run driven through the Python
mirror of the C-ABI (same config, same drand48 distortion) ending in the
same state."""
import ctypes
import os
import subprocess
import xml.etree.ElementTree as ET

import numpy as np
import pytest

import shell_amd as S
from conftest import plate

pytestmark = pytest.mark.gpu

H = 0.01
PCR = 2 * H * H / np.sqrt(3 * (1 - 0.3 ** 2))


def say(*a):
    print(*a, flush=True)


CONFIG = f"""# explicit run of a level-3 sphere under a pressure ramp, with self-contact
shell_mass_density   = 1
shell_young_modulus  = 1
shell_young_modulus2 = 0.5
shell_poisson_ratio  = 0.3
shell_shear_modulus  = 0.35
thickness            = {H}
damping_alpha        = 0.05
pressure             = '{float(-2 * PCR)!r}*min(t/4,1)'
shell_shell_contact  = linear
shell_cavity_contact = false
mu_s                 = 0.3
mu_d                 = 0.3
cell_size_mult       = 1
newmark_beta         = 0
newmark_gamma        = 0.5
dt                   = 0.05
adapt_dt_every       = 0
time_steps           = 200
data_time            = 1.0
vtk_time             = 2.0
state_time           = 0
random_seed          = 3
distortion           = 1e-2
"""


def _setup(tmp_path):
    X, T = S.mesh_icosphere(3)
    (tmp_path / "sphere.off").write_text(S.write_off(X, T))
    (tmp_path / "config.cfg").write_text(CONFIG)
    return X, T


def _run(tmp_path, *args):
    say("case driver:", " ".join(args))
    r = subprocess.run([S.DRIVER_PATH, "-m", "sphere.off", *args], cwd=tmp_path, capture_output=True, text=True,
                       timeout=120)
    return r


def _read_state(path):
    with open(path) as f:
        step = int(f.readline())
        dt, t, _ = (float(f.readline()) for _ in range(3))
        rows = []
        for line in f:
            vals = line.split()
            if len(vals) == 12:
                rows.append([float(x) for x in vals])
            else:
                break
    a = np.array(rows)
    u = a[:, [0, 3, 6]].reshape(-1)
    v = a[:, [1, 4, 7]].reshape(-1)
    return step, t, u, v


def test_driver_run_outputs_and_restart(tmp_path):
    _setup(tmp_path)
    r = _run(tmp_path)
    say(r.stdout[-1500:])
    assert r.returncode == 0, r.stdout + r.stderr
    # the cadence of shell.C:340-356: an output after the step whose time
    # t (accumulated step by step, t += dt) passes the next multiple of the
    # period, (unsigned)(t / period) growing
    t, data_steps, vtk_steps, ld, lv = 0.0, [], [0], 0, 0
    for k in range(1, 201):
        t += 0.05
        if int(t / 1.0) > ld:
            ld = int(t / 1.0)
            data_steps.append(k)
        if int(t / 2.0) > lv:
            lv = int(t / 2.0)
            vtk_steps.append(k)
    lines = [l for l in r.stdout.splitlines() if l.startswith("# ")]
    assert [int(l[2:].split()[1]) for l in lines] == [0] + data_steps
    cols = lines[-1][2:].split()
    assert len(cols) == 19 and float(cols[3]) == pytest.approx(10.0, rel=1e-12)
    data = tmp_path / "data"
    vtus = sorted(p.name for p in data.glob("shell*.vtu"))
    assert vtus == [f"shell{s:010d}.vtu" for s in vtk_steps]
    root = ET.parse(data / vtus[-1]).getroot()
    piece = root.find("UnstructuredGrid/Piece")
    assert piece.get("NumberOfPoints") == "642" and piece.get("NumberOfCells") == "1280"
    names = [a.get("Name") for a in piece.iter("DataArray")]
    assert names == ["ux", "uy", "uz", "ecav", "efric", "ekin", "eself", "mass", "ebend", "etens", "kGauss", "kMean",
                     "Points", "connectivity", "offsets", "types"]
    kg = np.array(piece.find("CellData/DataArray[@Name='kGauss']").text.split(), dtype=float)
    assert kg.size == 1280 and np.all(np.isfinite(kg))
    assert (data / "shell_state0000000200.dat").exists() and os.path.islink(data / "shell_state_latest.dat")
    step, t, u_full, _ = _read_state(data / "shell_state0000000200.dat")
    assert step == 200
    # the same run in two halves, the second one loading the first one's state file
    r1 = _run(tmp_path, "-d", "half", "-time_steps", "100")
    assert r1.returncode == 0, r1.stdout
    r2 = _run(tmp_path, "-d", "half2", "-time_steps", "200", "-l", "half/shell_state0000000100.dat")
    assert r2.returncode == 0, r2.stdout
    step2, t2, u_half, _ = _read_state(tmp_path / "half2" / "shell_state0000000200.dat")
    rel = np.linalg.norm(u_half - u_full) / np.linalg.norm(u_full)
    say("  restart: step", step2, "t", t2, "rel diff", rel)
    assert step2 == 200 and t2 == t
    assert rel < 1e-12


def test_driver_equals_the_python_mirror(tmp_path):
    """the driver's run equals the same run through ShellSystem: config read
    by the library, distortion u += distortion (2 drand48() - 1) h per DOF
    after srand48(random_seed) (ShellSystem.C:342-343, 863-869)"""
    X, T = _setup(tmp_path)
    r = _run(tmp_path, "-time_steps", "120", "-vtk_time", "0", "-data_time", "0")
    assert r.returncode == 0, r.stdout
    _, t, u_drv, v_drv = _read_state(tmp_path / "data" / "shell_state0000000120.dat")
    p, d, _ = S.read_config(CONFIG)
    libc = ctypes.CDLL("libc.so.6")
    libc.drand48.restype = ctypes.c_double
    libc.srand48(d["random_seed"])
    u0 = np.array([d["distortion"] * (2 * libc.drand48() - 1) * p["thickness"] for _ in range(3 * len(X))])
    g = S.ShellSystem(X, T, **p)
    g.set_state(u0)
    g.apply_initial_conditions()
    g.advance(120)
    u, v, _, tg = g.get_state()
    g.close()
    say("  driver vs mirror: t", t, tg, "max |du|", np.max(np.abs(u - u_drv)))
    assert tg == t
    assert np.array_equal(u, u_drv) and np.array_equal(v, v_drv)


CAVITY_CONFIG = """# a shell of radius 0.099 crumpled by a shrinking ellipsoidal cavity, in the
# reference's units and with its config.cfg keys (SI: Pa, kg/m^3, m, s)
shell_mass_density      = 1e3
shell_young_modulus     = 1e9
shell_young_modulus2    = 1e9
shell_poisson_ratio     = 1/3
shell_shear_modulus     = 3.75e8
thickness               = 1e-3
scale                   = 0.099,0.099,0.099
cavity                  = ellipsoid
Rx                      = 0.1*exp(-4*t)
Ry                      = 0.1*exp(-4*t)
Rz                      = 0.1*exp(-4*t)
shell_cavity_contact    = true
mu_s_cavity             = 0.1
mu_d_cavity             = 0.1
shell_shell_contact     = linear
damping_alpha           = 160
newmark_beta            = 0.25
dt                      = 1e-6
adapt_dt_every          = 1
rel_loc_err_max         = 5e-4
rel_loc_err_min         = 5e-5
friction_stop_velocity  = 1e-3
cell_margin             = 0.02
time_steps              = 600
data_time               = 0.0005
vtk_time                = 0
state_time              = 0
random_seed             = 0
distortion              = 1e-4
"""


def test_driver_cavity_run_equals_the_python_mirror(tmp_path):
    """the driver on a cavity configuration in the reference's units (the
    cavity closes on the shell after about 200 steps): property lines with
    cavity columns, the same state as the Python mirror's run, and cavity
    springs in the state file"""
    X, T = S.mesh_icosphere(3)
    (tmp_path / "sphere.off").write_text(S.write_off(X, T))
    (tmp_path / "config.cfg").write_text(CAVITY_CONFIG)
    say("case driver: cavity config, 600 steps")
    r = subprocess.run([S.DRIVER_PATH, "-m", "sphere.off"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = [l for l in r.stdout.splitlines() if l.startswith("# ")]
    cols = lines[-1][2:].split()
    say("  last property line:", lines[-1])
    # columns: date, step, dt, time, secs, area, volume, density, p_cav, Nc_cav, Nc, ...
    assert float(cols[7]) > 0 and float(cols[8]) > 0 and float(cols[9]) > 0
    _, t, u_drv, v_drv = _read_state(tmp_path / "data" / "shell_state0000000600.dat")
    p, d, _ = S.read_config(CAVITY_CONFIG)
    Xs = S.transform(X, scale=d["scale"])
    libc = ctypes.CDLL("libc.so.6")
    libc.drand48.restype = ctypes.c_double
    libc.srand48(d["random_seed"])
    u0 = np.array([d["distortion"] * (2 * libc.drand48() - 1) * p["thickness"] for _ in range(3 * len(X))])
    g = S.ShellSystem(Xs, T, **p)
    g.set_state(u0)
    g.apply_initial_conditions()
    g.advance(600)
    u, v, _, tg = g.get_state()
    springs = g.cavity_springs()
    g.close()
    say("  driver vs mirror: t", t, tg, "max |du|", np.max(np.abs(u - u_drv)), "springs", int(np.any(springs, 1).sum()))
    assert tg == t and np.array_equal(u, u_drv) and np.array_equal(v, v_drv)
    rows = np.array([[float(x) for x in l.split()] for l in
                     open(tmp_path / "data" / "shell_state0000000600.dat").read().splitlines()[4:4 + len(X)]])
    assert np.array_equal(rows[:, 9:12], springs)


PLATE_CONFIG = """# a square plate pinned on its edges, sagging under its own weight
shell_mass_density   = 1
shell_young_modulus  = 1
shell_young_modulus2 = 1
shell_poisson_ratio  = 0.3
shell_shear_modulus  = 1/2.6
thickness            = 0.05
gravity              = 3e-3
gravity_dir          = 1
damping_alpha        = 0.1
shell_boundary       = pinned
bc_every             = 1
shell_shell_contact  = none
shell_cavity_contact = false
newmark_beta         = 0
newmark_gamma        = 0.5
dt                   = 0.02
adapt_dt_every       = 0
time_steps           = 80
data_time            = 0
vtk_time             = 1.6
state_time           = 0
random_seed          = 5
distortion           = 1e-2
"""


def test_driver_open_mesh_equals_the_python_mirror(tmp_path):
    """an open mesh through the driver: its ghost layer added (shell.C:
    228-230), the pinned edges' conditions, the distortion over every node
    (ghosts included, ShellSystem.C:863-869), the .vtu files holding the
    real nodes only (ShellSystem.C:3189); equal to the same run through
    ShellSystem on the mesh with its ghost layer"""
    X, T = plate(8)
    (tmp_path / "sphere.off").write_text(S.write_off(X, T))
    (tmp_path / "config.cfg").write_text(PLATE_CONFIG)
    r = _run(tmp_path)
    assert r.returncode == 0, r.stdout
    _, t, u_drv, v_drv = _read_state(tmp_path / "data" / "shell_state0000000080.dat")
    Xo, To = S.read_off((tmp_path / "sphere.off").read_text())
    Xg, Tg, ng = S.mesh_add_ghosts(Xo, To)
    p, d, _ = S.read_config(PLATE_CONFIG)
    p["ghost_elems"] = ng
    libc = ctypes.CDLL("libc.so.6")
    libc.drand48.restype = ctypes.c_double
    libc.srand48(d["random_seed"])
    u0 = np.array([d["distortion"] * (2 * libc.drand48() - 1) * p["thickness"] for _ in range(3 * len(Xg))])
    g = S.ShellSystem(Xg, Tg, **p)
    g.set_state(u0)
    g.apply_initial_conditions()
    g.advance(80)
    u, v, _, tg = g.get_state()
    g.close()
    say("  open mesh driver vs mirror: t", t, tg, "max |du|", np.max(np.abs(u - u_drv)), "ghost triangles", ng)
    assert len(u_drv) == 3 * len(Xg) and tg == t
    assert np.array_equal(u, u_drv) and np.array_equal(v, v_drv)
    vtu = sorted((tmp_path / "data").glob("shell*.vtu"))
    assert vtu
    piece = ET.parse(vtu[-1]).getroot().find("UnstructuredGrid/Piece")
    assert int(piece.get("NumberOfPoints")) == len(X) and int(piece.get("NumberOfCells")) == len(T)
