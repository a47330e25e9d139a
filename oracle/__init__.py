"""CPU FP64 oracle of the explicit force-and-step path (TEST INFRASTRUCTURE).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this package, and only as the checker or as the timed CPU baseline.
The product library never loads it.

Each oracle function is pinned by tests/test_oracle_pins.py (closed forms,
invariants, brute force, and the SHA-256 of the reference's mesh files);
tests/test_golden.py keeps it reproducing the committed fixtures.  The
displacement rebuild trigger (an extension) is pinned by brute force: every
pair within h is listed at every checkpoint, and the trajectory equals the
reference trigger's.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def load(path):
    """a handle on an oracle library at `path` (the tests load copies built
    with a planted slip through this, to show that a pin catches it)"""
    L = ctypes.CDLL(path)
    _declare(L)
    return L


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


class Params(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in (
        "E1", "E2", "nu", "G12", "thickness", "rho", "memb_stiff_mult", "bend_stiff_mult",
        "damping_alpha", "damping_beta", "damping_gamma", "pressure_final", "pressure_ramp_time",
        "bulk_modulus", "volume_final_ratio", "volume_ramp_time", "contact_stiffness", "mu_s", "mu_d",
        "self_damping", "self_friction_damping", "fric_stiff_mult", "friction_stop_velocity",
        "static_friction_tol", "cell_size_mult", "cell_margin", "neighbor_margin", "newmark_beta",
        "newmark_gamma", "dt", "dt_min", "dt_max", "rel_loc_err_min", "rel_loc_err_max")] + [
        (n, ctypes.c_int32) for n in ("adapt_dt_every",
        "contact", "quirk_fulltime", "quirk_volume_schedule", "omp_faithful", "volume_at_prediction",
        "rebuild_trigger", "cavity_contact", "cavity", "gravity_dir", "pad0")] + [
        ("cavity_R0", ctypes.c_double * 3)] + [(n, ctypes.c_double) for n in (
        "cavity_rate", "mu_s_cavity", "mu_d_cavity", "cavity_damping", "cavity_friction_damping", "gravity",
        "damping_rotation")] + [(n, ctypes.c_int32) for n in ("growth_base", "thickness_dir", "grow_mass", "pad1")] + [
        (n, ctypes.c_double * 2) for n in ("growth_g0", "growth_gt", "growth_gx", "growth_gy", "growth_gz")] + [
        ("metric_growth_rate", ctypes.c_double), ("curvature_growth_rate", ctypes.c_double)] + [
        (n, ctypes.c_int32) for n in ("ghost_elems", "shell_boundary", "bc_coords", "bc_every")] + [
        ("bc_stiff_mult", ctypes.c_double)]


class Report(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in (
        "t", "residual", "fnorm", "fext_norm", "e_tens", "e_bend", "e_self", "e_fric", "e_kin",
        "volume", "area", "du_max")] + [(n, ctypes.c_int32) for n in (
        "n_contacts", "n_pairs", "n_rebuilds", "problem", "n_rejected")] + [("dt", ctypes.c_double),
                                                                          ("residual_deform", ctypes.c_double),
                                                                          ("e_cav", ctypes.c_double),
                                                                          ("n_cav", ctypes.c_int32),
                                                                          ("pad", ctypes.c_int32)]


_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int32)
_U = ctypes.POINTER(ctypes.c_uint32)
_V = ctypes.c_void_p


def _declare(L):
    L.orc_create.restype = _V
    L.orc_create.argtypes = [ctypes.c_int, ctypes.c_int, _D, _I, ctypes.POINTER(Params)]
    L.orc_destroy.argtypes = [_V]
    L.orc_boundary_edges.argtypes = [_V, _I]
    L.orc_bc_keep.argtypes = [_V, ctypes.POINTER(ctypes.c_int8)]
    L.orc_last_error.restype = ctypes.c_char_p
    for name in ("orc_patch_size",):
        getattr(L, name).argtypes = [_V, _I]
    L.orc_patches.argtypes = [_V, _I, _I]
    L.orc_elem_tables.argtypes = [_V, _D]
    L.orc_mass.argtypes = [_V, _D, _D]
    L.orc_set_state.argtypes = [_V, _D, _D, _D, ctypes.c_double]
    L.orc_get_state.argtypes = [_V, _D, _D, _D, _D]
    L.orc_forces.argtypes = [_V, _D, _D, ctypes.c_double, _D, _D, _D, _D]
    L.orc_internal_force.argtypes = [_V, _D, _D]
    L.orc_init_dynamics.argtypes = [_V]
    L.orc_step.argtypes = [_V, ctypes.c_int, ctypes.POINTER(Report)]
    L.orc_measure.argtypes = [_V]
    L.orc_properties.argtypes = [_V, _D]
    L.orc_node_cavity.argtypes = [_V, _D, _D, _D]
    L.orc_elem_growth.argtypes = [_V, _D]
    L.orc_elem_measure.argtypes = [_V, _D]
    L.orc_contact_rebuild.argtypes = [_V, ctypes.c_int]
    L.orc_pairs.argtypes = [_V, _U, _U, ctypes.c_int]
    L.orc_pair_state.argtypes = [_V, _D, _I, _D, ctypes.c_int]
    L.orc_pair_coords.argtypes = [_V, _D, ctypes.c_int]
    L.orc_smooth_contact.argtypes = [_V, ctypes.c_int, ctypes.c_int, _D, _D]
    L.orc_limit_point.argtypes = [ctypes.c_int, _D, ctypes.c_double, ctypes.c_double, _D]
    L.orc_limit_point.restype = None
    L.orc_smooth_closest.argtypes = [ctypes.c_int, _D, ctypes.c_int, _D, _D, _D]
    L.orc_smooth_closest.restype = None
    L.orc_close_counts.argtypes = [_V, _I]
    L.orc_close_sets.argtypes = [_V, _I, _I, ctypes.c_int]
    L.orc_cfl_omega_max.restype = ctypes.c_double
    L.orc_cfl_omega_max.argtypes = [_V, ctypes.c_int, ctypes.c_uint64]
    L.orc_icosphere_counts.argtypes = [ctypes.c_int, _I, _I]
    L.orc_icosphere.argtypes = [ctypes.c_int, ctypes.c_int, _D, _I]
    L.orc_stam_basis.argtypes = [ctypes.c_double, ctypes.c_double, _D]
    L.orc_raw_elem_basis.argtypes = [ctypes.c_int, _D]
    L.orc_loop_subdiv_rows.argtypes = [ctypes.c_int, _D]
    L.orc_tri_distance.argtypes = [_D, _D]
    L.orc_segment_distance.argtypes = [_D, _D]
    L.orc_point_triangle.argtypes = [_D, _D]
    L.orc_friction_apply.argtypes = [_D]
    L.orc_contact_force.restype = ctypes.c_double
    L.orc_contact_force.argtypes = [ctypes.c_double] * 3
    L.orc_contact_energy.restype = ctypes.c_double
    L.orc_contact_energy.argtypes = [ctypes.c_double] * 3


def dp(a):
    return a.ctypes.data_as(_D)


def ip(a):
    return a.ctypes.data_as(_I)


def up_(a):
    return a.ctypes.data_as(_U)


def icosphere(level, round_g=True):
    nv = ctypes.c_int32()
    ne = ctypes.c_int32()
    lib().orc_icosphere_counts(level, ctypes.byref(nv), ctypes.byref(ne))
    X = np.zeros((nv.value, 3))
    T = np.zeros((ne.value, 3), dtype=np.int32)
    lib().orc_icosphere(level, 1 if round_g else 0, dp(X), ip(T))
    return X, T


def make_params(**kw):
    p = Params()
    defaults = default_params()
    defaults.update(kw)
    for k, v in defaults.items():
        if k == "cavity_R0":
            v = (ctypes.c_double * 3)(*v)
        elif k.startswith("growth_g"):
            v = (ctypes.c_double * 2)(*v)
        setattr(p, k, v)
    return p


def default_params():
    return dict(E1=1.0, E2=0.5, nu=0.3, G12=0.35, thickness=0.01, rho=1.0, memb_stiff_mult=1.0,
                bend_stiff_mult=1.0, damping_alpha=0.05, damping_beta=0.0, damping_gamma=0.0,
                pressure_final=0.0, pressure_ramp_time=0.0, bulk_modulus=0.0, volume_final_ratio=1.0,
                volume_ramp_time=0.0, contact_stiffness=0.0, mu_s=0.0, mu_d=0.0, self_damping=0.0,
                self_friction_damping=0.0, fric_stiff_mult=1.0, friction_stop_velocity=1e-3,
                static_friction_tol=0.1, cell_size_mult=4.0, cell_margin=0.02, neighbor_margin=0.4,
                newmark_beta=0.0, newmark_gamma=0.5, dt=1e-3, dt_min=0.0, dt_max=0.0, rel_loc_err_min=5e-5,
                rel_loc_err_max=5e-4, adapt_dt_every=0, contact=0, quirk_fulltime=0,
                quirk_volume_schedule=0, omp_faithful=0, volume_at_prediction=0,
                rebuild_trigger=0, cavity_contact=0, cavity=2, gravity_dir=2, pad0=0,
                cavity_R0=(1.0, 1.0, 1.0), cavity_rate=0.0, mu_s_cavity=0.0, mu_d_cavity=0.0,
                cavity_damping=0.0, cavity_friction_damping=0.0, gravity=0.0, damping_rotation=0.0,
                growth_base=0, thickness_dir=2, grow_mass=1, pad1=0, growth_g0=(1.0, 1.0), growth_gt=(0.0, 0.0),
                growth_gx=(0.0, 0.0), growth_gy=(0.0, 0.0), growth_gz=(0.0, 0.0), metric_growth_rate=0.0,
                curvature_growth_rate=0.0, ghost_elems=0, shell_boundary=4, bc_coords=7, bc_every=0,
                bc_stiff_mult=1.0)


class Oracle:
    """Handle on one oracle shell system (FP64, CPU)."""

    def __init__(self, X, T, lib_handle=None, **params):
        self.L = lib_handle if lib_handle is not None else lib()
        self.X = np.ascontiguousarray(X, dtype=np.float64)
        self.T = np.ascontiguousarray(T, dtype=np.int32)
        self.params = make_params(**params)
        self.nv, self.ne = len(self.X), len(self.T) - self.params.ghost_elems    # ne: the real triangles
        self.h = self.L.orc_create(self.nv, len(self.T), dp(self.X), ip(self.T), ctypes.byref(self.params))
        if not self.h:
            raise RuntimeError("oracle: " + self.L.orc_last_error().decode())

    def close(self):
        if self.h:
            self.L.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def boundary_edges(self):
        """(m, 4) BoundaryCache nodes n1 n2 n3 n4 of the boundary edges, in order"""
        m = self.L.orc_boundary_edges(self.h, None)
        out = np.zeros((m, 4), dtype=np.int32)
        self.L.orc_boundary_edges(self.h, ip(out))
        return out

    def bc_keep(self, keep):
        """keep the boundary edges whose flag is set (the bc_criterion filter)"""
        k = np.ascontiguousarray(keep, dtype=np.int8)
        self.L.orc_bc_keep(self.h, k.ctypes.data_as(ctypes.POINTER(ctypes.c_int8)))

    def patches(self):
        idx = np.zeros((self.ne, 12), dtype=np.int32)
        tri = np.zeros((self.ne, 3), dtype=np.int32)
        self.L.orc_patches(self.h, ip(idx), ip(tri))
        N = np.zeros(self.ne, dtype=np.int32)
        self.L.orc_patch_size(self.h, ip(N))
        return idx, tri, N

    def elem_tables(self):
        out = np.zeros((self.ne, 101))
        self.L.orc_elem_tables(self.h, dp(out))
        return out

    def mass(self):
        Minv = np.zeros(3 * self.nv)
        C = np.zeros(3 * self.nv)
        self.L.orc_mass(self.h, dp(Minv), dp(C))
        return Minv, C

    def set_state(self, u, v=None, a=None, t=0.0):
        z = np.zeros(3 * self.nv)
        u = np.ascontiguousarray(u, dtype=np.float64).reshape(-1)
        v = z if v is None else np.ascontiguousarray(v, dtype=np.float64).reshape(-1)
        a = z if a is None else np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
        self.L.orc_set_state(self.h, dp(u), dp(v), dp(a), t)

    def get_state(self):
        u, v, a = (np.zeros(3 * self.nv) for _ in range(3))
        t = ctypes.c_double()
        self.L.orc_get_state(self.h, dp(u), dp(v), dp(a), ctypes.byref(t))
        return u, v, a, t.value

    def forces(self, u_pred, v=None, t=0.0):
        u_pred = np.ascontiguousarray(u_pred, dtype=np.float64).reshape(-1)
        vv = None if v is None else np.ascontiguousarray(v, dtype=np.float64).reshape(-1)
        f = np.zeros(3 * self.nv)
        fext = np.zeros(3 * self.nv)
        en = np.zeros(4)
        area = np.zeros(self.ne)
        self.L.orc_forces(self.h, dp(u_pred), dp(vv) if vv is not None else None, t, dp(f), dp(fext),
                         dp(en), dp(area))
        return f, fext, en, area

    def internal_force(self, u_pred):
        u_pred = np.ascontiguousarray(u_pred, dtype=np.float64).reshape(-1)
        f = np.zeros(3 * self.nv)
        self.L.orc_internal_force(self.h, dp(u_pred), dp(f))
        return f

    def init_dynamics(self):
        self.L.orc_init_dynamics(self.h)

    def step(self, n=1):
        rep = Report()
        self.L.orc_step(self.h, n, ctypes.byref(rep))
        return {k: getattr(rep, k) for k, _ in Report._fields_}

    def measure(self):
        self.L.orc_measure(self.h)

    def properties(self):
        out = np.zeros(16)
        self.L.orc_properties(self.h, dp(out))
        return out

    def elem_growth(self):
        """per element G (3), GV (3 x 3 row-major), alpha (3), beta (3)"""
        out = np.zeros((self.ne, 18))
        self.L.orc_elem_growth(self.h, dp(out))
        return out[:, :3], out[:, 3:12].reshape(-1, 3, 3), out[:, 12:15], out[:, 15:18]

    def node_cavity(self):
        """per node: cavity spring origin (nv x 3), cavity and cavity-friction
        energy over the node's mean element area, from the last evaluation"""
        o = np.zeros((self.nv, 3))
        ec = np.zeros(self.nv)
        ef = np.zeros(self.nv)
        self.L.orc_node_cavity(self.h, dp(o), dp(ec), dp(ef))
        return o, ec, ef

    def elem_measure(self):
        out = np.zeros((self.ne, 10))
        self.L.orc_elem_measure(self.h, dp(out))
        return out

    def rebuild(self, force=True):
        return self.L.orc_contact_rebuild(self.h, 1 if force else 0)

    def pairs(self):
        n = self.L.orc_pairs(self.h, None, None, 0)
        e1 = np.zeros(max(n, 1), dtype=np.uint32)
        e2 = np.zeros(max(n, 1), dtype=np.uint32)
        self.L.orc_pairs(self.h, up_(e1), up_(e2), n)
        return e1[:n], e2[:n]

    def pair_state(self):
        n = self.L.orc_pairs(self.h, None, None, 0)
        d2 = np.zeros(max(n, 1))
        st = np.zeros(max(n, 1), dtype=np.int32)
        so = np.zeros((max(n, 1), 3))
        self.L.orc_pair_state(self.h, dp(d2), ip(st), dp(so), n)
        return d2[:n], st[:n], so[:n]

    def pair_coords(self):
        """smooth contact: v1 w1 v2 w2 of each pair's last closest points (n x 4)"""
        n = self.L.orc_pairs(self.h, None, None, 0)
        vw = np.zeros((max(n, 1), 4))
        self.L.orc_pair_coords(self.h, dp(vw), n)
        return vw[:n]

    def smooth_contact(self, e1, e2, x=(1 / 3, 1 / 3, 1 / 3, 1 / 3)):
        """SmoothShellShellContact(e1, e2) + compute() from x at the state's
        positions: dict of d2, dc, point, x (the minimiser), iterations"""
        xs = np.array(x, dtype=np.float64)
        out = np.zeros(8)
        if self.L.orc_smooth_contact(self.h, e1, e2, dp(xs), dp(out)) != 0:
            raise ValueError((e1, e2))
        return {"d2": out[0], "dc": out[1:4], "point": out[4:7], "x": xs, "iterations": int(out[7])}

    def close_sets(self):
        start = np.zeros(self.ne + 1, dtype=np.int32)
        n = self.L.orc_close_sets(self.h, ip(start), None, 0)
        ids = np.zeros(max(n, 1), dtype=np.int32)
        self.L.orc_close_sets(self.h, ip(start), ip(ids), n)
        return start, ids[:n]

    def cfl_omega_max(self, iters=60, seed=1):
        return self.L.orc_cfl_omega_max(self.h, iters, seed)


def stam_basis(v, w):
    out = np.zeros(72)
    lib().orc_stam_basis(v, w, dp(out))
    return out.reshape(6, 12)


def raw_elem_basis(valence):
    out = np.zeros(6 * (valence + 6))
    if lib().orc_raw_elem_basis(valence, dp(out)) != 0:
        raise ValueError(valence)
    return out.reshape(6, valence + 6)


def subdiv_rows(valence):
    S = np.zeros(12 * (valence + 6))
    lib().orc_loop_subdiv_rows(valence, dp(S))
    return S.reshape(12, valence + 6)


def limit_point(valence, X, xi, eta, handle=None):
    """the oracle's limit surface of a patch (points X, (valence+6) x 3) at
    (xi, eta): 6 x 3 = point, d/dxi, d/deta, xixi, etaeta, xieta"""
    X = np.ascontiguousarray(X, dtype=np.float64).reshape(-1)
    out = np.zeros(18)
    (handle or lib()).orc_limit_point(valence, dp(X), xi, eta, dp(out))
    return out.reshape(6, 3)


def smooth_closest(valence1, X1, valence2, X2, x=(1 / 3, 1 / 3, 1 / 3, 1 / 3), handle=None):
    """the oracle's closest points of two limit-surface patches: dict of d2,
    dc, x (the minimiser), iterations"""
    X1 = np.ascontiguousarray(X1, dtype=np.float64).reshape(-1)
    X2 = np.ascontiguousarray(X2, dtype=np.float64).reshape(-1)
    xs = np.array(x, dtype=np.float64)
    out = np.zeros(5)
    (handle or lib()).orc_smooth_closest(valence1, dp(X1), valence2, dp(X2), dp(xs), dp(out))
    return {"d2": out[0], "dc": out[1:4], "x": xs, "iterations": int(out[4])}


def tri_distance(P):
    P = np.ascontiguousarray(P, dtype=np.float64).reshape(18)
    out = np.zeros(8)
    lib().orc_tri_distance(dp(P), dp(out))
    return out


def segment_distance(P):
    P = np.ascontiguousarray(P, dtype=np.float64).reshape(12)
    out = np.zeros(5)
    lib().orc_segment_distance(dp(P), dp(out))
    return out


def point_triangle(P):
    P = np.ascontiguousarray(P, dtype=np.float64).reshape(12)
    out = np.zeros(5)
    lib().orc_point_triangle(dp(P), dp(out))
    return out


def friction_apply(io):
    io = np.ascontiguousarray(io, dtype=np.float64).copy()
    lib().orc_friction_apply(dp(io))
    return io


def contact_force(k, h, d):
    return lib().orc_contact_force(k, h, d)


def contact_energy(k, h, d):
    return lib().orc_contact_energy(k, h, d)
